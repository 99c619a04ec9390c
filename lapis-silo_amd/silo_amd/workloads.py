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
