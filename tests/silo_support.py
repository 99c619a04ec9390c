"""The reference's saved database state (testBaseData/siloSerializedState, written by
database.test.cpp:71-96 from testBaseData/unitTestDummyDataset) and the same five rows built from
that dataset's input.ndjson with the oracle's diff, for the loader tests."""
import datetime
import gzip
import json
import os
import shutil

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATE = os.path.join(GOLDEN, "silo_state")
DATASET = os.path.join(GOLDEN, "unit_test_dataset")
TIMESTAMP = "1785915539"


def saved_state(tmp_path) -> str:
    """A copy of the saved state under tmp_path (the schema archive is kept gzip-compressed)."""
    root = os.path.join(str(tmp_path), "silo_state")
    d = os.path.join(root, TIMESTAMP)
    os.makedirs(d, exist_ok=True)
    src = os.path.join(STATE, TIMESTAMP)
    for name in os.listdir(src):
        if name.endswith(".gz"):
            with gzip.open(os.path.join(src, name), "rb") as f, open(os.path.join(d, name[:-3]), "wb") as g:
                shutil.copyfileobj(f, g)
        else:
            shutil.copy(os.path.join(src, name), os.path.join(d, name))
    return d


def rows() -> list:
    with open(os.path.join(DATASET, "input.ndjson")) as f:
        return [json.loads(line) for line in f if line.strip()]


def references() -> dict:
    with open(os.path.join(DATASET, "reference_genomes.json")) as f:
        g = json.load(f)
    out = {s["name"]: (0, s["sequence"]) for s in g["nucleotideSequences"]}
    out.update({s["name"]: (1, s["sequence"]) for s in g["genes"]})
    return out


def days(date: str | None) -> int:
    if date is None:
        return 0
    return (datetime.date.fromisoformat(date) - datetime.date(1970, 1, 1)).days


def seq_column(name: str, alphabet: int, reference: str, sequences: list):
    """A sequence column in diff form from aligned strings (None = null), diffed by the oracle."""
    from oracle import oracle
    from silo_amd import table

    n = len(sequences)
    is_null = np.zeros(n, np.uint8)
    cs = np.zeros(n, np.uint32)
    ce = np.zeros(n, np.uint32)
    mptr = np.zeros(n + 1, np.uint64)
    nptr = np.zeros(n + 1, np.uint64)
    mpos, msym, ns, ne = [], [], [], []
    for i, s in enumerate(sequences):
        if s is None:
            is_null[i] = 1
        else:
            start, end, missing, muts = oracle.diff_sequence(s, reference, alphabet)
            cs[i], ce[i] = start, end
            for p, sym in muts:
                mpos.append(p)
                msym.append(sym)
            for p in missing:
                if ne and ne[-1] == p and len(ns) > nptr[i]:
                    ne[-1] = p + 1
                else:
                    ns.append(p)
                    ne.append(p + 1)
        mptr[i + 1] = len(mpos)
        nptr[i + 1] = len(ns)
    ref = np.array([table.symbol_of(alphabet, c) for c in reference], np.uint8)
    return table.SeqColumnData(name, alphabet, ref, is_null, cs, ce, mptr, np.array(mpos, np.uint32),
                               np.array(msym, np.uint8), nptr, np.array(ns, np.uint32), np.array(ne, np.uint32))


def lineage_definition(text: str) -> tuple[list, np.ndarray, np.ndarray]:
    """Names, single parents (-1 = root or several parents) and alias targets (-1 = none) from a
    lineage definition file's YAML (lineage name -> {parents: [...], aliases: [...]})."""
    import yaml

    doc = yaml.safe_load(text)
    names = list(doc)
    aliases = [(a, k) for k, v in doc.items() for a in ((v or {}).get("aliases") or [])]
    names += [a for a, _ in aliases]
    index = {n: i for i, n in enumerate(names)}
    parent = np.full(len(names), -1, np.int32)
    for k, v in doc.items():
        ps = (v or {}).get("parents") or []
        if len(ps) == 1:
            parent[index[k]] = index[ps[0]]
    alias = np.full(len(names), -1, np.int32)
    for a, k in aliases:
        alias[index[a]] = index[k]
    return names, parent, alias


def ndjson_spec(lineage_text: str, aliases: bool = True):
    """The dataset's rows as a table in diff form: the four sequence columns, date, unsorted_date and
    age, and the dictionary columns (pango_lineage with the lineage definition's tree, and its
    aliases unless `aliases` is false)."""
    from silo_amd import table

    rs = rows()
    n = len(rs)
    spec = table.TableSpec(chunk_sizes=np.asarray([n], np.uint32))
    for name, (alphabet, ref) in references().items():
        spec.seq_columns.append(seq_column(name, alphabet, ref, [(r.get(name) or {}).get("sequence") for r in rs]))
    for name, is_date in (("date", True), ("unsorted_date", True), ("age", False)):
        vals = [r.get(name) for r in rs]
        nulls = np.array([v is None for v in vals], np.uint8)
        ints = np.array([days(v) if is_date else (v or 0) for v in vals], np.int32)
        spec.int_columns.append(table.IntColumnData(name, is_date, ints, nulls))
    for name in ("country", "division", "region"):
        vals = [r.get(name) for r in rs]
        names = sorted({v for v in vals if v is not None})
        ids = np.array([names.index(v) if v is not None else 0 for v in vals], np.uint32)
        spec.dict_columns.append(table.DictColumnData(name, ids, np.array([v is None for v in vals], np.uint8), names))
    names, parent, alias = lineage_definition(lineage_text)
    vals = [r.get("pango_lineage") for r in rs]
    ids = np.array([names.index(v) for v in vals], np.uint32)
    spec.dict_columns.append(table.DictColumnData("pango_lineage", ids, np.zeros(n, np.uint8), names, parent,
                                                  alias if aliases else None))
    return spec
