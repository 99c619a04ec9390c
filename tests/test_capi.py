"""The C-ABI library loads and exports every entry point include/rdb.h declares; the host-only
helpers behave (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rdb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rdb_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from silo_amd import _native

    lib = ctypes.CDLL(_native.ENGINE_LIB)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from silo_amd.engine import SIGNATURES

    assert sorted(n for n, _, _ in SIGNATURES) == declared_symbols()


def test_engine_diff_matches_oracle_diff():
    from oracle.oracle import diff_sequence as oracle_diff
    from silo_amd.engine import diff_sequence

    rng = np.random.default_rng(3)
    ref = "".join(rng.choice(list("ACGT"), 200))
    for _ in range(200):
        s = list(ref)
        for _ in range(rng.integers(0, 10)):
            s[rng.integers(0, 200)] = rng.choice(list("ACGTN-RYacgtnu"))
        if rng.random() < 0.3:
            a = rng.integers(0, 200)
            s[a:a + rng.integers(1, 40)] = ["N"] * len(s[a:a + 40])
        s = "".join(s)[:200]
        assert diff_sequence(s, ref) == oracle_diff(s, ref)
    assert diff_sequence("NNNN", "ACGT") == (0, 0, [], [])


def test_engine_diff_rejects_illegal_character():
    from silo_amd.engine import EngineError, diff_sequence

    with pytest.raises(EngineError) as e:
        diff_sequence("AC!T", "ACGT")
    assert e.value.status == "RDB_EINVAL"
