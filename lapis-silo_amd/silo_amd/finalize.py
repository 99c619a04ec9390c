"""The finalized state of a sequence column, computed on the host from its diff form.

`SequenceColumn::finalize` (`sequence_column.cpp:157-212`) turns the appended rows into what a saved
state holds: the local reference adapted position by position (a symbol replaces the current one
when strictly more covered rows carry it, `vertical_sequence_index.cpp:56-164`), the vertical index
of every (position, chunk, symbol) other than the local reference with its rows, run-optimized, and
per row the covered range and the missing positions inside it
(`horizontal_coverage_index.cpp:17-46`).  This module computes that state with numpy so a table in
diff form can be written as a saved state (silo_archive.save_database); the engine's own index
build runs on the GPU.
"""
from __future__ import annotations

import numpy as np

from .table import ALPHABET_NUC, FinalizedSeqColumnData, SeqColumnData


def _rows_of(ptr: np.ndarray) -> np.ndarray:
    counts = np.diff(np.asarray(ptr, np.int64))
    return np.repeat(np.arange(len(counts), dtype=np.int64), counts)


def _container(rows: np.ndarray) -> tuple[int, bytes]:
    """(typecode, container_write bytes) of sorted in-chunk rows: an array up to 4,096 rows, else a
    bitset, and runs where those are smaller (CRoaring's runOptimize)."""
    card = len(rows)
    size = 2 * card if card <= 4096 else 8192
    r = rows.astype(np.int64)
    breaks = np.nonzero(np.diff(r) != 1)[0]
    starts = np.concatenate([r[:1], r[breaks + 1]])
    ends = np.concatenate([r[breaks], r[-1:]])
    if 2 + 4 * len(starts) < size:
        body = np.stack([starts, ends - starts], 1).astype("<u2").tobytes()
        return 3, np.uint16(len(starts)).astype("<u2").tobytes() + body
    if card <= 4096:
        return 2, r.astype("<u2").tobytes()
    bits = np.zeros(65536, np.uint8)
    bits[r] = 1
    return 1, np.packbits(bits, bitorder="little").tobytes()


def finalize_column(col: SeqColumnData, chunk_sizes: np.ndarray) -> FinalizedSeqColumnData:
    chunk_sizes = np.asarray(chunk_sizes, np.int64)
    n = int(chunk_sizes.sum())
    L = len(col.reference)
    nsym = 16 if col.alphabet == ALPHABET_NUC else 28
    ref = np.asarray(col.reference, np.int64)
    null = np.asarray(col.is_null).astype(bool)
    start = np.where(null, 0, np.asarray(col.cov_start, np.int64))
    end = np.where(null, 0, np.asarray(col.cov_end, np.int64))
    offsets = np.concatenate([[0], np.cumsum(chunk_sizes)])
    chunk_of = np.repeat(np.arange(len(chunk_sizes), dtype=np.int64), chunk_sizes)
    in_chunk = np.arange(n, dtype=np.int64) - offsets[chunk_of]

    # missing runs trimmed to the covered range (insertCoverage's removeRange), non-null rows only
    run_row = _rows_of(col.nrun_ptr)
    ra = np.maximum(np.asarray(col.nrun_start, np.int64), start[run_row])
    rb = np.minimum(np.asarray(col.nrun_end, np.int64), end[run_row])
    keep = (ra < rb) & ~null[run_row]
    run_row, ra, rb = run_row[keep], ra[keep], rb[keep]

    # coverage per position (computeCoverageCardinalities)
    count = lambda x: np.bincount(x, minlength=L + 1).astype(np.int64)
    coverage = np.cumsum(count(start) - count(end) - count(ra) + count(rb))[:L]

    mut_row = _rows_of(col.mut_ptr)
    mpos = np.asarray(col.mut_pos, np.int32)
    msym = np.asarray(col.mut_sym, np.int32)
    if null.any():
        live = ~null[mut_row]
        mut_row, mpos, msym = mut_row[live], mpos[live], msym[live]
    counts = np.bincount(mpos * nsym + msym, minlength=L * nsym).reshape(L, nsym).astype(np.int64)
    stored = counts.sum(1)
    cur = counts.copy()
    cur[np.arange(L), ref] = coverage - stored + counts[np.arange(L), ref]
    # the first symbol (other than the current one) with the highest count replaces the current one
    # if it has strictly more rows
    others = cur.copy()
    others[np.arange(L), ref] = -1
    best = others.argmax(1)
    flip = others[np.arange(L), best] > cur[np.arange(L), ref]
    local = np.where(flip, best, ref).astype(np.uint8)

    # entries (position, chunk, symbol, row in chunk): stored mutations except those to the new local
    # reference, and the original reference's rows at flipped positions
    keep = ~(flip[mpos] & (msym == best[mpos]))
    e_pos, e_sym, e_row = mpos[keep], msym[keep], mut_row[keep]
    extra = []
    for p in np.nonzero(flip)[0].tolist():
        covered = (start <= p) & (p < end)
        missing_rows = run_row[(ra <= p) & (p < rb)]
        covered[missing_rows] = False
        covered[mut_row[mpos == p]] = False
        rows = np.nonzero(covered)[0]
        extra.append((np.full(len(rows), p, np.int64), np.full(len(rows), int(ref[p]), np.int64), rows))
    if extra:
        e_pos = np.concatenate([e_pos] + [x[0] for x in extra])
        e_sym = np.concatenate([e_sym] + [x[1] for x in extra])
        e_row = np.concatenate([e_row] + [x[2] for x in extra])
    # one 64-bit key per entry: position (16 bits) | chunk (16) | symbol (8) | row in chunk (16)
    full = (e_pos.astype(np.uint64) << np.uint64(40)) | (chunk_of[e_row].astype(np.uint64) << np.uint64(24)) | \
        (e_sym.astype(np.uint64) << np.uint64(16)) | in_chunk[e_row].astype(np.uint64)
    del e_pos, e_sym, e_row
    full.sort()
    key = (full >> np.uint64(16)).astype(np.int64)
    e_low = (full & np.uint64(0xFFFF)).astype(np.int64)
    del full
    e_pos = key >> 24
    e_chunk = (key >> 8) & 0xFFFF
    e_sym = key & 0xFF
    cuts = np.nonzero(np.diff(key))[0] + 1
    # container boundaries in the sorted entries (no entries: no containers)
    bounds = np.concatenate([[0], cuts, [len(key)]]).astype(np.int64) if len(key) else np.zeros(1, np.int64)
    k = len(bounds) - 1
    first = bounds[:-1]
    card = np.diff(bounds)
    # container types as runOptimize leaves them: runs where 2 + 4 runs bytes beat the array (2 per
    # row, up to 4,096 rows) or bitset (8 KiB) form
    owner = np.repeat(np.arange(k, dtype=np.int64), card)
    run_head = np.ones(len(e_low), bool)
    run_head[1:] = (owner[1:] != owner[:-1]) | (e_low[1:] != e_low[:-1] + 1)
    n_runs = np.bincount(owner[run_head], minlength=k)
    plain_size = np.where(card <= 4096, 2 * card, 8192)
    is_run = 2 + 4 * n_runs < plain_size
    tc = np.where(is_run, 3, np.where(card <= 4096, 2, 1)).astype(np.uint8)
    size = np.where(is_run, 2 + 4 * n_runs, plain_size)
    offs = np.zeros(k, np.int64)
    offs[1:] = np.cumsum(size)[:-1]
    payload = np.zeros(int(size.sum()), np.uint8)
    words = payload.view("<u2")  # every container size is even
    # consecutive array containers are consecutive rows in both the entries and the payload
    prev = 0
    for i in np.nonzero(tc != 2)[0].tolist() + [k]:
        if prev < i:
            a0, a1 = int(first[prev]), int(bounds[i])
            w0 = int(offs[prev]) // 2
            words[w0:w0 + a1 - a0] = e_low[a0:a1]
        if i < k:
            t, body = _container(e_low[bounds[i]:bounds[i + 1]])
            assert t == tc[i]
            payload[offs[i]:offs[i] + len(body)] = np.frombuffer(body, np.uint8)
        prev = i + 1
    card = card.astype(np.uint32)
    offs = offs.astype(np.uint64)

    # missing runs per row, ascending, touching runs joined (a roaring bitmap holds them as one)
    order = np.lexsort((ra, run_row))
    run_row, ra, rb = run_row[order], ra[order], rb[order]
    head = np.ones(len(ra), bool)
    head[1:] = (run_row[1:] != run_row[:-1]) | (ra[1:] > rb[:-1])
    heads = np.nonzero(head)[0]
    run_row, ra, rb = run_row[heads], ra[heads], np.maximum.reduceat(rb, heads) if len(heads) else rb[:0]
    rptr = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(run_row, minlength=n), out=rptr[1:])

    return FinalizedSeqColumnData(
        col.name, col.alphabet, np.asarray(col.reference, np.uint8), local, e_pos[first].astype(np.uint32),
        e_chunk[first].astype(np.uint16), e_sym[first].astype(np.uint8), tc, card, offs, payload,
        null.astype(np.uint8), start.astype(np.uint32),
        end.astype(np.uint32), None, None, rptr, ra.astype(np.uint32), rb.astype(np.uint32))
