"""The query workloads of the benchmark configurations.

* config 0 (100 k rows): Mutations(minProportion 0.05) under And(date in a 90-day range,
  country == most frequent value), plus Mutations under single nucleotide-equals and
  has-mutation filters.
* config 1 (4 M rows): 64 Mutations(minProportion 0.05) queries whose filters are date ranges,
  country equality or lineage equality, with target selectivities log-spaced over 1e-4 .. 0.5.
  The filter kind cycles date -> country -> lineage; each picks the parameter whose expected
  selectivity (from the generator's distributions) is closest to the target, and a date range is
  narrowed by a country equality when one day is still too broad.
"""
from __future__ import annotations

import math

import numpy as np

from .datagen import Dataset

MIN_PROPORTION = 0.05


def config0_queries(ds: Dataset, n_positions: int = 6, seed: int = 0) -> list[dict]:
    rng = np.random.default_rng(seed)
    top = ds.country_names[int(np.argmax(ds.country_weight))]
    d0 = int(ds.spec.date_base + ds.spec.n_days // 3)
    out = [{"filter": f'(and (date_between date {d0} {d0 + 89}) (str_eq country "{top}"))', "kind": "date90&country"}]
    positions = rng.choice(np.unique(ds.mut_pos), size=n_positions, replace=False)
    for p in positions.tolist():
        sym = "ACGT-"[int(rng.integers(0, 5))]
        out.append({"filter": f"(sym_eq main {p} {sym})", "kind": "nucleotideEquals"})
        out.append({"filter": f"(has_mut main {p})", "kind": "hasMutation"})
    for q in out:
        q["min_proportion"] = MIN_PROPORTION
    return out


def config1_queries(ds: Dataset, n: int = 64) -> list[dict]:
    cw = ds.country_weight / ds.country_weight.sum()
    lw = ds.lineage_weight / ds.lineage_weight.sum()
    n_days = ds.spec.n_days
    base = ds.spec.date_base
    targets = np.logspace(math.log10(1e-4), math.log10(0.5), n)
    out = []
    for i, s in enumerate(targets.tolist()):
        kind = ("date", "country", "lineage")[i % 3]
        if kind == "country" and s < cw.min() / 2:
            kind = "lineage"
        if kind == "date":
            days = s * n_days
            start = base + (i * 37) % max(n_days - 1, 1)
            if days >= 1:
                w = int(round(days))
                start = min(start, base + n_days - w)
                f = f"(date_between date {start} {start + w - 1})"
                expect = w / n_days
            else:
                # one day narrowed by the country whose share best matches
                c = int(np.argmin(np.abs(np.log(cw * (1.0 / n_days)) - math.log(s))))
                f = f'(and (date_between date {start} {start}) (str_eq country "{ds.country_names[c]}"))'
                expect = cw[c] / n_days
        elif kind == "country":
            c = int(np.argmin(np.abs(np.log(cw) - math.log(s))))
            f = f'(str_eq country "{ds.country_names[c]}")'
            expect = cw[c]
        else:
            v = int(np.argmin(np.abs(np.log(lw) - math.log(s))))
            f = f'(lineage lineage "{ds.lineage_names[v]}" 0)'
            expect = lw[v]
        out.append({"filter": f, "kind": kind, "target_selectivity": float(s), "expected_selectivity": float(expect),
                    "min_proportion": MIN_PROPORTION})
    return out


class TreeGen:
    """Random filter trees of the batched workload (config 2): depth 1-6 over And / Or / Not /
    NOf, n-of-m with m uniform in 2..20 and n uniform in 1..m; leaves are NucleotideSymbolEquals
    (symbol from A, C, G, T, -, N) or HasMutation at uniform positions, or metadata predicates
    (country or lineage equality, date range)."""

    SYMBOLS = "ACGT-N"

    def __init__(self, ds: Dataset, seed: int):
        self.rng = np.random.default_rng(seed)
        self.ds = ds
        # positions drawn uniformly, but half of the sequence leaves at positions some row mutates,
        # so that leaves are not almost always trivially empty or full
        self.hot = np.unique(ds.mut_pos)

    def position(self) -> int:
        r = self.rng
        if len(self.hot) and r.random() < 0.5:
            return int(self.hot[int(r.integers(0, len(self.hot)))])
        return int(r.integers(0, self.ds.genome_length))

    def leaf(self) -> str:
        r = self.rng
        k = r.random()
        if k < 0.35:
            return f"(sym_eq main {self.position()} {self.SYMBOLS[int(r.integers(0, 6))]})"
        if k < 0.6:
            return f"(has_mut main {self.position()})"
        if k < 0.75:
            c = int(r.integers(0, len(self.ds.country_names)))
            return f'(str_eq country "{self.ds.country_names[c]}")'
        if k < 0.85:
            v = int(r.integers(0, len(self.ds.lineage_names)))
            return f'(lineage lineage "{self.ds.lineage_names[v]}" {int(r.integers(0, 2))})'
        a = self.ds.spec.date_base + int(r.integers(0, self.ds.spec.n_days))
        return f"(date_between date {a} {a + int(r.integers(0, 200))})"

    def tree(self, depth: int) -> str:
        r = self.rng
        if depth <= 1:
            return self.leaf()
        k = r.random()
        if k < 0.35:
            parts = [self.tree(depth - 1)] + [self.tree(int(r.integers(1, depth))) for _ in range(int(r.integers(1, 3)))]
            return "(and " + " ".join(parts) + ")"
        if k < 0.6:
            parts = [self.tree(depth - 1)] + [self.tree(int(r.integers(1, depth))) for _ in range(int(r.integers(1, 3)))]
            return "(or " + " ".join(parts) + ")"
        if k < 0.75:
            return f"(not {self.tree(depth - 1)})"
        m = int(r.integers(2, 21))
        n = int(r.integers(1, m + 1))
        # children: mostly leaves, one subtree carries the depth
        parts = [self.tree(depth - 1)] + [self.tree(int(r.integers(1, min(depth, 3)))) for _ in range(m - 1)]
        return f"(nof {n} {int(r.random() < 0.2)} " + " ".join(parts) + ")"


def config2_queries(ds: Dataset, count, n: int = 1024, seed: int = 2, lo: float = 1e-6, hi: float = 0.5,
                    candidates: int = 12000, min_proportion: float = MIN_PROPORTION) -> list[dict]:
    """1,024 heterogeneous filter trees whose exact selectivities fill log-uniform bins over
    [lo, hi]: candidates are drawn from TreeGen, their exact cardinalities come from `count` (a
    function from a list of filters to their cardinalities, e.g. the engine's batched Count), and
    each of n equal-width log10 bins takes its share of the candidates that fall in it (bins left
    short are topped up from their neighbours).  Queries alternate between Mutations
    (minProportion 0.05) and Aggregated by country; each also returns its cardinality."""
    gen = TreeGen(ds, seed)
    filters = [gen.tree(int(gen.rng.integers(1, 7))) for _ in range(candidates)]
    cards = np.asarray(count(filters), dtype=np.float64)
    sel = cards / ds.n_rows
    n_bins = 32
    edges = np.linspace(math.log10(lo), math.log10(hi), n_bins + 1)
    ok = (sel >= lo) & (sel <= hi)
    bins = np.clip(np.searchsorted(edges, np.log10(np.where(ok, sel, 1.0)), side="right") - 1, 0, n_bins - 1)
    per_bin = [list(np.flatnonzero(ok & (bins == b))) for b in range(n_bins)]
    rng = np.random.default_rng(seed + 1)
    for lst in per_bin:
        rng.shuffle(lst)
    want = [n // n_bins + (1 if b < n % n_bins else 0) for b in range(n_bins)]
    chosen = []
    spare = []
    for b in range(n_bins):
        take = per_bin[b][: want[b]]
        chosen += take
        spare += per_bin[b][want[b]:]
    rng.shuffle(spare)
    chosen += spare[: n - len(chosen)]
    chosen = chosen[:n]
    out = []
    for i, idx in enumerate(chosen):
        q = {"filter": filters[idx], "selectivity": float(sel[idx]), "cardinality": int(cards[idx])}
        if i % 2 == 0:
            q.update(action="mutations", column="main", min_proportion=min_proportion)
        else:
            q.update(action="aggregate", dims=["country"])
        out.append(q)
    return out
