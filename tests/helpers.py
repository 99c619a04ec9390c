"""Shared test helpers: small seeded datasets in the engine's ingest form, the same data for the
oracle and the brute-force checker, and a seeded generator of random filter expressions."""
from __future__ import annotations

import random

import numpy as np

from silo_amd import datagen, table

NUC = "-ACGTRYSWKMBDHVN"


def small_spec(seed: int = 7, n_rows: int = 3000, L: int = 400, chunk: int = 1024, lineage: bool = True,
               null_seq: float = 0.02, null_meta: float = 0.05):
    spec = datagen.GenSpec(n_rows=n_rows, genome_length=L, seed=seed, lineage_model=lineage, n_nodes=30, n_mean=20,
                           null_seq_frac=null_seq, null_meta_frac=null_meta, n_days=50)
    ds = datagen.generate(spec, keep_handle=True)
    ts = table.spec_from_dataset(ds, chunk=chunk)
    return ds, ts


def brute_for(ds, ts):
    from oracle import brute

    rows = [None if ds.is_null[i] else ds.materialize(i) for i in range(ds.n_rows)]
    ref = "".join(NUC[x] for x in ds.reference)
    return brute.BruteTable(
        ts.chunk_sizes, seq_columns={"main": (0, ref, rows)},
        int_columns={"date": (ds.date, ds.date_null)},
        dict_columns={"country": (ds.country, ds.country_null, ds.country_names, None),
                      "lineage": (ds.lineage, ds.lineage_null, ds.lineage_names, ds.lineage_parent)},
        bitmaps=ts.bitmaps)


class ExprGen:
    """Random filter expressions over the columns of `spec_from_dataset` tables."""

    def __init__(self, ds, seed: int, bitmaps: int = 0):
        self.r = random.Random(seed)
        self.ds = ds
        self.L = ds.genome_length
        self.bitmaps = bitmaps
        self.date_lo = int(ds.date[ds.date_null == 0].min()) if ds.n_rows else 0
        self.date_hi = int(ds.date[ds.date_null == 0].max()) if ds.n_rows else 0
        # positions that carry mutations are more interesting than random ones
        self.hot = sorted(set(ds.mut_pos[: 5000].tolist())) or list(range(self.L))

    def pos(self) -> int:
        return self.r.choice(self.hot) if self.r.random() < 0.7 else self.r.randrange(self.L)

    def leaf(self) -> str:
        r = self.r
        k = r.randrange(10 if self.bitmaps else 9)
        if k == 0:
            return f"(sym_eq main {self.pos()} {r.choice('ACGT-N.')})"
        if k == 1:
            return f"(has_mut main {self.pos()})"
        if k == 2:
            chars = "".join(sorted(set(r.choice("ACGT-NRY") for _ in range(r.randrange(1, 4)))))
            return f'(sym_in main {self.pos()} "{chars}")'
        if k == 3:
            v = r.choice(self.ds.country_names + ["null", "nowhere"])
            return "(str_eq country null)" if v == "null" else f'(str_eq country "{v}")'
        if k == 4:
            v = r.choice(self.ds.lineage_names)
            return f'(lineage lineage "{v}" {r.randrange(2)})'
        if k == 5:
            a = r.randrange(self.date_lo, self.date_hi + 1)
            b = a + r.randrange(0, 20)
            lo = "_" if r.random() < 0.1 else str(a)
            hi = "_" if r.random() < 0.1 else str(b)
            return f"(date_between date {lo} {hi})"
        if k == 6:
            return r.choice(["(true)", "(false)"])
        if k == 7:
            return f"(covered main {self.pos()})"
        if k == 8:
            return f"(sym_eq main {self.pos()} {r.choice('ACGT')})"
        return f"(bitmap {r.randrange(self.bitmaps)})"

    def expr(self, depth: int = 3) -> str:
        r = self.r
        if depth == 0 or r.random() < 0.3:
            return self.leaf()
        k = r.randrange(4)
        if k == 0:
            return "(and " + " ".join(self.expr(depth - 1) for _ in range(r.randrange(1, 4))) + ")"
        if k == 1:
            return "(or " + " ".join(self.expr(depth - 1) for _ in range(r.randrange(1, 4))) + ")"
        if k == 2:
            return f"(not {self.expr(depth - 1)})"
        m = r.randrange(1, 6)
        n = r.randrange(-1, m + 2)
        return f"(nof {n} {r.randrange(2)} " + " ".join(self.expr(depth - 1) for _ in range(m)) + ")"


def random_bitmaps(ts, seed: int, count: int = 3) -> list:
    rng = np.random.default_rng(seed)
    ids = ts.global_ids()
    out = []
    for _ in range(count):
        k = int(rng.integers(0, max(len(ids) // 3, 1)))
        out.append(np.sort(rng.choice(ids, size=k, replace=False)).astype(np.uint32))
    return out


def column_from_matrix(name: str, reference: np.ndarray, M: np.ndarray, null: np.ndarray, alphabet: int = 0):
    """The diff form (aligned_sequence.cpp:18-122 with offset 0) of aligned rows given as a symbol
    matrix M[row, position]: leading and trailing missing symbols bound the covered range, missing
    symbols inside it become runs, other symbols that differ from the reference are mutations.
    Null rows (null[i] != 0) get the empty range [0, 0).  Vectorised, for tables of 10^5 rows."""
    missing = 15 if alphabet == 0 else 27
    n, L = M.shape
    isn = (M == missing) | null.astype(bool)[:, None]
    covered_any = ~isn.all(axis=1)
    start = np.where(covered_any, np.argmax(~isn, axis=1), 0).astype(np.uint32)
    end = np.where(covered_any, L - np.argmax(~isn[:, ::-1], axis=1), 0).astype(np.uint32)
    col = np.arange(L, dtype=np.uint32)[None, :]
    inside = (col >= start[:, None]) & (col < end[:, None])
    nmask = isn & inside
    prev = np.zeros_like(nmask)
    prev[:, 1:] = nmask[:, :-1]
    nxt = np.zeros_like(nmask)
    nxt[:, :-1] = nmask[:, 1:]
    rs, cs = np.nonzero(nmask & ~prev)
    re_, ce = np.nonzero(nmask & ~nxt)
    assert np.array_equal(rs, re_)
    nrun_ptr = np.zeros(n + 1, np.uint64)
    nrun_ptr[1:] = np.cumsum(np.bincount(rs, minlength=n))
    mm = inside & ~isn & (M != reference[None, :])
    mr, mc = np.nonzero(mm)
    mut_ptr = np.zeros(n + 1, np.uint64)
    mut_ptr[1:] = np.cumsum(np.bincount(mr, minlength=n))
    return table.SeqColumnData(name, alphabet, reference.astype(np.uint8), null.astype(np.uint8), start, end, mut_ptr,
                               mc.astype(np.uint32), M[mr, mc].astype(np.uint8), nrun_ptr, cs.astype(np.uint32),
                               (ce + 1).astype(np.uint32))


def bitset_table(seed: int = 5):
    """Three chunks (65,536 + 65,536 + 40,000 rows) x 96 positions whose vertical index holds bitset
    containers (more than 4,096 scattered rows of a non-local-reference symbol in one chunk), arrays
    just under 4,096 rows, run containers, a position whose local reference flips, internal missing
    runs, ambiguity codes and null rows; a 3-value dictionary column whose value bitmaps are
    bitsets; a date column.  Returns (spec, M, null) with M the symbol matrix."""
    rng = np.random.default_rng(seed)
    sizes = np.asarray([65536, 65536, 40000], np.uint32)
    n, L = int(sizes.sum()), 96
    ref = np.tile(np.asarray([1, 2, 3, 4], np.uint8), L // 4)  # ACGT...
    M = np.repeat(ref[None, :], n, axis=0)
    r = np.arange(n)
    local = np.concatenate([np.arange(s) for s in sizes.tolist()])
    A, C, G, T, DEL, R, N = 1, 2, 3, 4, 0, 5, 15
    M[r % 8 == 3, 10] = T                               # 8,192 rows per full chunk, 8,192 runs: bitset
    M[rng.random(n) < 0.3, 20] = C                      # ~19.7 k random rows per chunk: bitset
    M[r % 16 == 0, 30] = A                              # 4,096 rows per full chunk less nulls: array
    M[r % 16 == 1, 30] = T
    M[(r % 16 == 2) | (r % 128 == 7), 31] = DEL         # 4,608 rows per full chunk less nulls: bitset
    M[((local >= 1000) & (local < 9000)) | ((local >= 30000) & (local < 31000)), 40] = G  # runs
    u = rng.random(n)
    M[u < 0.42, 50] = A                                 # A is the majority: G (the reference) is stored
    M[(u >= 0.42) & (u < 0.8), 50] = C
    M[r % 50 == 1, 5] = R                               # ambiguity code
    M[(r % 3 == 0) & (r % 13 != 0), 62] = C             # bitset next to missing runs
    M[r % 13 == 0, 60:64] = N                           # internal missing runs
    M[r % 97 == 5, 0:3] = N                             # leading missing symbols
    M[r % 89 == 7, 94:96] = N                           # trailing missing symbols
    for lo, hi in ((70, 82), (82, 94)):                 # two private substitutions per row
        p = rng.integers(lo, hi, n)
        M[r, p] = np.asarray([1, 2, 3, 4, 0], np.uint8)[(ref[p].astype(np.int64) + rng.integers(0, 4, n)) % 5]
    null = (r % 211 == 17).astype(np.uint8)
    M[null.astype(bool)] = N
    spec = table.TableSpec(chunk_sizes=sizes)
    spec.seq_columns.append(column_from_matrix("main", ref, M, null))
    ids = rng.integers(0, 3, n).astype(np.uint32)
    cnull = (rng.random(n) < 0.02).astype(np.uint8)
    spec.dict_columns.append(table.DictColumnData("country", ids, cnull, ["c0", "c1", "c2"]))
    dnull = (rng.random(n) < 0.03).astype(np.uint8)
    spec.int_columns.append(table.IntColumnData("date", True, (18262 + rng.integers(0, 100, n)).astype(np.int32),
                                                dnull))
    return spec, M, null


BITSET_FILTERS = ["(true)", "(sym_eq main 10 T)", "(sym_eq main 10 G)", "(sym_eq main 10 .)", "(has_mut main 20)",
                  "(sym_eq main 50 C)", "(sym_eq main 50 A)", '(sym_in main 50 "CG")', "(not (sym_eq main 31 -))",
                  "(sym_eq main 30 A)", "(or (sym_eq main 40 G) (sym_eq main 30 T))", "(has_mut main 5)",
                  "(sym_eq main 61 N)", "(covered main 62)", "(sym_eq main 62 C)", '(str_eq country "c1")',
                  '(and (str_eq country "c0") (has_mut main 62))', "(str_eq country null)",
                  "(nof 2 0 (has_mut main 10) (has_mut main 20) (has_mut main 50) (sym_eq main 62 C))",
                  "(nof 2 1 (sym_eq main 10 T) (not (sym_eq main 20 C)) (sym_eq main 31 -))",
                  "(date_between date 18270 18300)", "(and (has_mut main 31) (not (has_mut main 10)))"]


def brute_from_matrix(spec, M, null):
    from oracle import brute

    c = spec.seq_columns[0]
    ref = "".join(NUC[x] for x in c.reference)
    d = spec.dict_columns[0]
    i = spec.int_columns[0]
    return brute.BruteTable(spec.chunk_sizes, seq_columns={c.name: (0, ref, (M, null))},
                            int_columns={i.name: (i.values, i.is_null)},
                            dict_columns={d.name: (d.ids, d.is_null, d.names, None)})


def sorted_date_table(seed: int = 3):
    """A table whose `date` column is null-free and non-decreasing (the reference then compiles
    date ranges to RangeSelection, date_between.cpp:75-79): chunks of 3,000, 0, 2,500, 4,096 and
    1,500 rows, the values of chunk 0 all below the first value of chunk 2, repeated values across
    chunk boundaries; `date_u` holds the same values with one null (Selection path); `country` a
    small dictionary.  Returns (spec, date values)."""
    rng = np.random.default_rng(seed)
    sizes = np.asarray([3000, 0, 2500, 4096, 1500], np.uint32)
    n = int(sizes.sum())
    first = np.sort(rng.integers(18000, 18100, 3000))
    rest = np.sort(rng.integers(18100, 18400, n - 3000))
    rest[:40] = 18100                                    # a run of equal values at the start of chunk 2
    rest[2500 - 30:2500 + 30] = 18250                    # equal values across the chunk 2 / 3 boundary
    dates = np.sort(np.concatenate([first, rest])).astype(np.int32)
    spec = table.TableSpec(chunk_sizes=sizes)
    spec.int_columns.append(table.IntColumnData("date", True, dates, np.zeros(n, np.uint8)))
    u_null = np.zeros(n, np.uint8)
    u_null[4321] = 1
    spec.int_columns.append(table.IntColumnData("date_u", True, dates.copy(), u_null))
    spec.dict_columns.append(table.DictColumnData("country", rng.integers(0, 3, n).astype(np.uint32),
                                                  np.zeros(n, np.uint8), ["c0", "c1", "c2"]))
    return spec, dates


def sorted_date_filters(dates) -> list:
    """date_between filters whose ends fall below, at and above the column's values and exactly at
    chunk boundaries, open bounds, negations and combinations."""
    lo, hi = int(dates.min()), int(dates.max())
    edge = int(dates[2999])  # last value of chunk 0 (ranges ending exactly at its end)
    points = sorted({lo - 1, lo, lo + 1, edge, edge + 1, 18100, 18250, 18251, hi - 1, hi, hi + 1})
    out = []
    for a in points:
        out += [f"(date_between date {a} _)", f"(date_between date _ {a})", f"(not (date_between date {a} _))"]
        for b in points:
            if b >= a:
                out.append(f"(date_between date {a} {b})")
    out += ["(date_between date _ _)", f"(date_between date {hi + 5} {hi + 9})", f"(date_between date {lo} {lo - 3})",
            f'(and (date_between date 18100 18250) (str_eq country "c1"))',
            f"(or (date_between date _ {edge}) (date_between date 18300 _))",
            f"(not (or (date_between date _ 18050) (date_between date 18250 18260)))",
            f"(nof 2 0 (date_between date 18050 18300) (date_between date_u 18100 _) (str_eq country \"c0\"))",
            f"(date_between date_u 18100 18250)", f"(not (date_between date_u 18100 18250))"]
    return out
