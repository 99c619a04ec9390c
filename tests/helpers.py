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
