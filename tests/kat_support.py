"""Loads the reference known-answer vectors (tests/golden/reference_kats.json) into TableSpecs."""
import json
import os

import numpy as np

from silo_amd import table

KATS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")))


def dataset_spec(name: str) -> table.TableSpec:
    d = KATS["datasets"][name]
    spec = table.TableSpec(chunk_sizes=np.asarray(d["chunks"], np.uint32))
    n = spec.n_rows
    for col, (alphabet, ref, rows) in d.get("seq", {}).items():
        spec.seq_columns.append(table.seq_column_from_strings(col, ref, rows, alphabet))
    for col, values in d.get("dict", {}).items():
        names = sorted(set(v for v in values if v is not None))
        ids = np.array([names.index(v) if v is not None else 0 for v in values], np.uint32)
        nulls = np.array([v is None for v in values], np.uint8)
        spec.dict_columns.append(table.DictColumnData(col, ids, nulls, names))
    for col, values in d.get("date", {}).items():
        spec.int_columns.append(table.IntColumnData(col, True, np.asarray(values, np.int32), np.zeros(n, np.uint8)))
    return spec


def operator_spec(case) -> table.TableSpec:
    spec = table.TableSpec(chunk_sizes=np.asarray(case["chunks"], np.uint32))
    spec.bitmaps = [np.asarray(b, np.uint32) for b in case["bitmaps"]]
    return spec


def row_ids(spec, indices):
    g = spec.global_ids()
    return [int(g[i]) for i in indices]


def symbol_chars(alphabet: int) -> str:
    return table.alphabet_chars(alphabet)
