"""ORACLE — test infrastructure only.  Brute-force semantics over materialized sequences.

This module pins the C++ oracle without sharing any of its structure: it evaluates every filter
expression row by row on the full aligned sequences (numpy), counts symbols per position directly,
and groups rows with a dictionary.  It restates only the *meaning* the reference gives each
expression and action (no indexes, no local reference, no operator trees):

* symbol of a row at position p: null row -> none; N (missing) when p is outside the covered range
  or an N inside it; otherwise the character (`aligned_sequence.cpp:18-122`,
  `symbol_in_set.cpp:95-226`);
* `hasMutation(p)`: symbol not in AMBIGUITY_SYMBOLS[reference[p]] (`has_mutation.cpp:34-67`);
* NOf: number of matching children >= n, or == n when exact (`nof.cpp`, `threshold.cpp`);
* Mutations: per position, counts of each symbol over non-null filter rows, `coverage` = sum over
  the valid mutation symbols, threshold `ceil(total * p) - 1` in double (`mutations_node.cpp:292-366`);
* Aggregated: group counts ordered by value string (null last) per dimension, depth first
  (`bitmap_aggregation_node.cpp:91-130, 226-253`).

Suitable for a few thousand rows.
"""
from __future__ import annotations

import math

import numpy as np

NUC = "-ACGTRYSWKMBDHVN"
AA = "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X"

# IUPAC: what each symbol stands for
_NUC_CODES = {"-": "-", "A": "A", "C": "C", "G": "G", "T": "T", "R": "AG", "Y": "CT", "S": "GC", "W": "AT",
              "K": "GT", "M": "AC", "B": "CGT", "D": "AGT", "H": "ACT", "V": "ACG", "N": NUC}
_AA_CODES = {c: c for c in AA}
_AA_CODES.update({"B": "DN", "J": "LI", "Z": "QE", "X": AA})


def ambiguity(alphabet: int, sym: int) -> set:
    chars = NUC if alphabet == 0 else AA
    codes = _NUC_CODES if alphabet == 0 else _AA_CODES
    mine = set(codes[chars[sym]])
    return {i for i, c in enumerate(chars) if mine <= set(codes[c])}


def to_symbol(alphabet: int, ch: str) -> int:
    c = ch.upper()
    if alphabet == 0 and c == "U":
        c = "T"
    return (NUC if alphabet == 0 else AA).index(c)


class BruteTable:
    def __init__(self, chunk_sizes, seq_columns=None, int_columns=None, dict_columns=None, bitmaps=None):
        """seq_columns: {name: (alphabet, reference_str, rows)} with rows either [str | None per row]
        or (symbol matrix [n_rows, L], null flags [n_rows]);
        int_columns: {name: (values, is_null)}; dict_columns: {name: (ids, is_null, names, parent)}"""
        self.chunk_sizes = [int(x) for x in chunk_sizes]
        self.n = sum(self.chunk_sizes)
        self.gids = np.concatenate([(c << 16) + np.arange(s, dtype=np.int64)
                                    for c, s in enumerate(self.chunk_sizes)]) if self.n else np.zeros(0, np.int64)
        self.seq = {}
        for name, (alphabet, ref, rows) in (seq_columns or {}).items():
            L = len(ref)
            missing = (15 if alphabet == 0 else 27)
            if isinstance(rows, tuple):
                null = np.asarray(rows[1]).astype(bool)
                M = np.where(null[:, None], missing, np.asarray(rows[0], np.uint8)).astype(np.uint8)
                self.seq[name] = (alphabet, np.array([to_symbol(alphabet, c) for c in ref], np.uint8), M, null)
                continue
            M = np.full((self.n, L), missing, dtype=np.uint8)
            null = np.zeros(self.n, bool)
            for i, s in enumerate(rows):
                if s is None:
                    null[i] = True
                    continue
                M[i] = [to_symbol(alphabet, ch) for ch in s]
            self.seq[name] = (alphabet, np.array([to_symbol(alphabet, c) for c in ref], np.uint8), M, null)
        self.ints = {k: (np.asarray(v[0]), np.asarray(v[1]).astype(bool)) for k, v in (int_columns or {}).items()}
        self.dicts = {}
        self.aliases = {}
        for k, col in (dict_columns or {}).items():
            ids, nulls, names, parent = col[:4]
            self.dicts[k] = (np.asarray(ids), np.asarray(nulls).astype(bool), list(names),
                             None if parent is None else [int(x) for x in parent])
            # optional fifth entry: per value the value its lineage alias resolves to (-1 = none)
            if len(col) > 4 and col[4] is not None:
                self.aliases[k] = [int(x) for x in col[4]]
        self.bitmaps = [set(int(x) for x in b) for b in (bitmaps or [])]

    # ---- expressions -----------------------------------------------------------------------
    def parse(self, text: str):
        toks = []
        i = 0
        while i < len(text):
            ch = text[i]
            if ch.isspace():
                i += 1
            elif ch in "()":
                toks.append(ch)
                i += 1
            elif ch == '"':
                j = text.index('"', i + 1)
                toks.append(("str", text[i + 1:j]))
                i = j + 1
            else:
                j = i
                while j < len(text) and not text[j].isspace() and text[j] not in "()":
                    j += 1
                toks.append(text[i:j])
                i = j
        pos = 0

        def expr():
            nonlocal pos
            assert toks[pos] == "("
            pos += 1
            op = toks[pos]
            pos += 1
            args = []
            while toks[pos] != ")":
                if toks[pos] == "(":
                    args.append(expr())
                else:
                    args.append(toks[pos])
                    pos += 1
            pos += 1
            return (op, args)

        return expr()

    def eval(self, text: str) -> np.ndarray:
        return self._eval(self.parse(text))

    @staticmethod
    def _atom(a):
        return a[1] if isinstance(a, tuple) else a

    def _eval(self, node) -> np.ndarray:
        op, args = node
        n = self.n
        if op == "true":
            return np.ones(n, bool)
        if op == "false":
            return np.zeros(n, bool)
        if op == "and":
            r = np.ones(n, bool)
            for a in args:
                r &= self._eval(a)
            return r
        if op == "or":
            r = np.zeros(n, bool)
            for a in args:
                r |= self._eval(a)
            return r
        if op == "not":
            return ~self._eval(args[0])
        if op == "nof":
            k, exact = int(args[0]), int(args[1]) != 0
            cnt = np.zeros(n, np.int64)
            for a in args[2:]:
                cnt += self._eval(a)
            return cnt == k if exact else cnt >= k
        if op in ("sym_eq", "sym_in", "has_mut", "covered"):
            col, p = self._atom(args[0]), int(args[1])
            alphabet, ref, M, null = self.seq[col]
            sym = M[:, p]
            missing = 15 if alphabet == 0 else 27
            if op == "covered":
                return (~null) & (sym != missing)
            if op == "has_mut":
                bad = ambiguity(alphabet, int(ref[p]))
                return (~null) & ~np.isin(sym, list(bad))
            if op == "sym_eq":
                ch = self._atom(args[2])
                target = {int(ref[p])} if ch == "." else {to_symbol(alphabet, ch)}
            else:
                target = {to_symbol(alphabet, c) for c in self._atom(args[2])}
            return (~null) & np.isin(sym, list(target))
        if op == "str_eq" or op == "lineage":
            column = self._atom(args[0])
            ids, nulls, names, parent = self.dicts[column]
            v = args[1]
            if not isinstance(v, tuple) and v == "null":
                return nulls.copy()
            name = self._atom(v)
            if name not in names:
                return np.zeros(n, bool)
            vid = names.index(name)
            if op == "lineage":
                # lineage_index.cpp:17-45: values and the query resolve their alias first
                alias = self.aliases.get(column)
                res = np.array([alias[w] if alias and alias[w] >= 0 else w for w in range(len(names))], np.int64)
                vid = int(res[vid])
                if int(args[2]) != 0:
                    def is_desc(w):
                        seen = 0
                        while w >= 0 and seen <= len(parent):
                            if w == vid:
                                return True
                            w = parent[w]
                            seen += 1
                        return False
                    ok = np.array([is_desc(int(res[w])) for w in range(len(names))], bool)
                    return (~nulls) & ok[ids]
                return (~nulls) & (res[ids] == vid)
            return (~nulls) & (ids == vid)
        if op in ("date_between", "int_between"):
            vals, nulls = self.ints[self._atom(args[0])]
            lo = None if args[1] == "_" else int(args[1])
            hi = None if args[2] == "_" else int(args[2])
            r = ~nulls
            if lo is not None:
                r &= vals >= lo
            if hi is not None:
                r &= vals <= hi
            return r
        if op == "int_eq":
            vals, nulls = self.ints[self._atom(args[0])]
            return (~nulls) & (vals == int(args[1]))
        if op == "bitmap":
            s = self.bitmaps[int(args[0])]
            return np.array([int(g) in s for g in self.gids], bool)
        raise ValueError(op)

    def filter(self, text: str) -> np.ndarray:
        return self.gids[self.eval(text)].astype(np.uint32)

    # ---- actions ---------------------------------------------------------------------------
    def mutation_counts(self, text: str, column: str) -> np.ndarray:
        alphabet, ref, M, null = self.seq[column]
        nsym = 16 if alphabet == 0 else 28
        missing = 15 if alphabet == 0 else 27
        rows = M[self.eval(text) & ~null]
        out = np.zeros((nsym, M.shape[1]), np.uint32)
        for s in range(nsym):
            if s != missing:
                out[s] = (rows == s).sum(axis=0)
        return out

    def mutations(self, text: str, column: str, min_proportion: float) -> list:
        alphabet, ref, M, null = self.seq[column]
        counts = self.mutation_counts(text, column)
        valid = [0, 1, 2, 3, 4] if alphabet == 0 else list(range(23)) + [26]
        out = []
        for p in range(M.shape[1]):
            total = int(sum(int(counts[s, p]) for s in valid))
            if total == 0:
                continue
            thr = 0 if min_proportion == 0 else int(math.ceil(float(total) * min_proportion) - 1)
            for s in valid:
                if s == int(ref[p]):
                    continue
                c = int(counts[s, p])
                if c > thr:
                    out.append((p + 1, int(ref[p]), s, c, total, c / total))
        return out

    def aggregate(self, text: str, dims: list) -> list:
        f = self.eval(text)
        keys = []
        orders = []
        for d in dims:
            if "@" in d:
                col, p = d.split("@")
                alphabet, ref, M, null = self.seq[col]
                k = [None if null[i] else int(M[i, int(p)]) for i in range(self.n)]
                keys.append(k)
                orders.append(lambda v: (1, 0) if v is None else (0, v))
            else:
                ids, nulls, names, parent = self.dicts[d]
                k = [None if nulls[i] else int(ids[i]) for i in range(self.n)]
                keys.append(k)
                orders.append(lambda v, names=names: (1, "") if v is None else (0, names[v]))
        counts = {}
        for i in np.nonzero(f)[0]:
            key = tuple(k[i] for k in keys)
            counts[key] = counts.get(key, 0) + 1
        return sorted(counts.items(), key=lambda kv: tuple(o(v) for o, v in zip(orders, kv[0])))
