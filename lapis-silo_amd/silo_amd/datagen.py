"""Seeded synthetic workloads: sequences in reference-diffed form plus metadata.

The rows are produced by `csrc/datagen.cpp` (one RNG stream per row) and returned as numpy arrays
in the ingest layout both the engine and the CPU oracle accept:

* `is_null[n]`, `cov_start[n]`, `cov_end[n]` — the covered range `[start, end)` of each row;
* `mut_ptr[n+1]`, `mut_pos[]`, `mut_sym[]` — mutations against the reference (CSR by row,
  positions strictly increasing, symbols in the reference's nucleotide order `-ACGTRYSWKMBDHVN`);
* `nrun_ptr[n+1]`, `nrun_start[]`, `nrun_end[]` — internal missing (N) runs `[start, end)`;
* metadata: `country` / `lineage` dictionary ids, `date` (days since 1970-01-01).

Rows are laid out in 2^16-row chunks (the reference's `RowId`, `storage/column/row_id.h:16-37`).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native

BASE_SEED = 0x51105EED
NUC_CHARS = "-ACGTRYSWKMBDHVN"
DATE_BASE = 18262  # 2020-01-01
CHUNK = 1 << 16


@dataclass(frozen=True)
class GenSpec:
    n_rows: int
    genome_length: int = 29903
    seed: int = BASE_SEED
    lineage_model: bool = False
    sub_lambda: float = 30.0
    n_nodes: int = 2000
    zipf_lineage: float = 1.2
    node_lambda: float = 3.0
    private_lambda: float = 2.0
    del_prob: float = 0.02
    del_max: int = 30
    n_lambda: float = 3.0
    n_mean: float = 200.0
    n_countries: int = 50
    zipf_country: float = 1.1
    n_days: int = 1000
    date_base: int = DATE_BASE
    null_seq_frac: float = 0.0
    null_meta_frac: float = 0.0


# The five workload configurations: 0 plain CPU-runnable, 1-3 the 4 M-row lineage data, 4 serving.
CONFIGS = {
    0: GenSpec(n_rows=100_000, seed=BASE_SEED + 0),
    1: GenSpec(n_rows=4_000_000, seed=BASE_SEED + 1, lineage_model=True),
    2: GenSpec(n_rows=4_000_000, seed=BASE_SEED + 1, lineage_model=True),
    3: GenSpec(n_rows=4_000_000, seed=BASE_SEED + 1, lineage_model=True),
    4: GenSpec(n_rows=16_000_000, seed=BASE_SEED + 4, lineage_model=True),
}


def chunk_sizes(n_rows: int, chunk: int = CHUNK) -> np.ndarray:
    full, rem = divmod(n_rows, chunk)
    sizes = [chunk] * full + ([rem] if rem else [])
    return np.asarray(sizes, dtype=np.uint32)


_SIGNATURES = False


def _lib() -> ctypes.CDLL:
    global _SIGNATURES
    lib = _native.load(_native.DATAGEN_LIB)
    if not _SIGNATURES:
        u32, u64, f64, i32, vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p
        lib.dg_generate.restype = vp
        lib.dg_generate.argtypes = [u32, u32, u64, ctypes.c_int, f64, u32, f64, f64, f64, f64, u32,
                                    f64, f64, u32, f64, u32, i32, f64, f64]
        lib.dg_free.argtypes = [vp]
        for name in ("dg_num_rows", "dg_genome_length", "dg_num_lineages"):
            getattr(lib, name).restype = u32
            getattr(lib, name).argtypes = [vp]
        for name in ("dg_num_mutations", "dg_num_nruns"):
            getattr(lib, name).restype = u64
            getattr(lib, name).argtypes = [vp]
        ptr_fns = ["dg_reference", "dg_is_null", "dg_cov_start", "dg_cov_end", "dg_mut_ptr", "dg_mut_pos",
                   "dg_mut_sym", "dg_nrun_ptr", "dg_nrun_start", "dg_nrun_end", "dg_country", "dg_date",
                   "dg_lineage", "dg_country_null", "dg_date_null", "dg_lineage_null", "dg_lineage_parent",
                   "dg_lineage_weight", "dg_country_weight", "dg_country_name_perm", "dg_lineage_name_perm"]
        for name in ptr_fns:
            getattr(lib, name).restype = vp
            getattr(lib, name).argtypes = [vp]
        lib.dg_materialize.restype = ctypes.c_int
        lib.dg_materialize.argtypes = [vp, u32, ctypes.c_char_p]
        lib.dg_render.restype = None
        lib.dg_render.argtypes = [u32, vp, ctypes.c_char_p, ctypes.c_char, u64, u64] + [vp] * 10
        lib.dg_translate.restype = ctypes.c_int
        lib.dg_translate.argtypes = [vp, u32, vp, vp]
        for name in ("dg_aa_reference", "dg_aa_cov_start", "dg_aa_cov_end", "dg_aa_mut_ptr", "dg_aa_mut_pos",
                     "dg_aa_mut_sym", "dg_aa_nrun_ptr", "dg_aa_nrun_start", "dg_aa_nrun_end"):
            getattr(lib, name).restype = vp
            getattr(lib, name).argtypes = [vp, u32]
        for name in ("dg_aa_num_mutations", "dg_aa_num_nruns"):
            getattr(lib, name).restype = u64
            getattr(lib, name).argtypes = [vp, u32]
        _SIGNATURES = True
    return lib


def _copy(ptr: int, n: int, dtype) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=dtype)
    nbytes = n * np.dtype(dtype).itemsize
    buf = (ctypes.c_char * nbytes).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype, count=n).copy()


@dataclass
class Dataset:
    spec: GenSpec
    reference: np.ndarray
    is_null: np.ndarray
    cov_start: np.ndarray
    cov_end: np.ndarray
    mut_ptr: np.ndarray
    mut_pos: np.ndarray
    mut_sym: np.ndarray
    nrun_ptr: np.ndarray
    nrun_start: np.ndarray
    nrun_end: np.ndarray
    country: np.ndarray
    country_null: np.ndarray
    date: np.ndarray
    date_null: np.ndarray
    lineage: np.ndarray
    lineage_null: np.ndarray
    lineage_parent: np.ndarray
    lineage_weight: np.ndarray
    country_weight: np.ndarray
    country_names: list = field(default_factory=list)
    lineage_names: list = field(default_factory=list)
    _handle: int | None = None

    @property
    def n_rows(self) -> int:
        return int(self.is_null.shape[0])

    @property
    def genome_length(self) -> int:
        return int(self.reference.shape[0])

    def chunk_sizes(self, chunk: int = CHUNK) -> np.ndarray:
        return chunk_sizes(self.n_rows, chunk)

    def materialize(self, row: int) -> str:
        """The aligned sequence string row `row` was diffed from (regenerated from its RNG stream)."""
        if self._handle is None:
            raise RuntimeError("dataset was released")
        buf = ctypes.create_string_buffer(self.genome_length)
        if _lib().dg_materialize(self._handle, row, buf) != 0:
            raise IndexError(row)
        return buf.raw.decode("ascii")

    def release(self) -> None:
        if self._handle is not None:
            _lib().dg_free(self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.release()
        except Exception:
            pass


# The ten amino-acid columns of the serving workload: names and residue counts (9,700 in total).
AA_GENES = [("ORF1a", 4405), ("ORF1b", 2728), ("S", 1273), ("ORF3a", 275), ("E", 75), ("M", 222), ("ORF6", 61),
            ("ORF7a", 121), ("ORF8", 121), ("N", 419)]


def gene_layout(genome_length: int, genes=AA_GENES, seed: int = BASE_SEED):
    """Seeded, non-overlapping codon start positions of the genes in genome order: the spare
    nucleotides (genome length minus three per residue) are cut at seeded points into gaps."""
    total = 3 * sum(n for _, n in genes)
    spare = genome_length - total
    if spare < 0:
        raise ValueError("genes do not fit the genome")
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.integers(0, spare + 1, len(genes)))
    starts = []
    pos = 0
    prev = 0
    for (name, n), c in zip(genes, cuts):
        pos += int(c - prev)
        prev = int(c)
        starts.append(pos)
        pos += 3 * n
    return [(name, s, n) for (name, n), s in zip(genes, starts)]


def translate(ds: "Dataset", layout) -> list:
    """Amino-acid columns of the dataset's rows (the generator keeps its handle for this): per gene
    (name, reference, cov_start, cov_end, mut_ptr, mut_pos, mut_sym, nrun_ptr, nrun_start, nrun_end)."""
    if ds._handle is None:
        raise RuntimeError("generate with keep_handle=True to translate")
    lib = _lib()
    starts = np.asarray([s for _, s, _ in layout], np.uint32)
    lens = np.asarray([n for _, _, n in layout], np.uint32)
    if lib.dg_translate(ds._handle, len(layout), starts.ctypes.data, lens.ctypes.data) != 0:
        raise ValueError("gene layout outside the genome")
    h = ds._handle
    n = ds.n_rows
    out = []
    for g, (name, _, gl) in enumerate(layout):
        nm = lib.dg_aa_num_mutations(h, g)
        nr = lib.dg_aa_num_nruns(h, g)
        out.append((name, _copy(lib.dg_aa_reference(h, g), gl, np.uint8), _copy(lib.dg_aa_cov_start(h, g), n, np.uint32),
                    _copy(lib.dg_aa_cov_end(h, g), n, np.uint32), _copy(lib.dg_aa_mut_ptr(h, g), n + 1, np.uint64),
                    _copy(lib.dg_aa_mut_pos(h, g), nm, np.uint32), _copy(lib.dg_aa_mut_sym(h, g), nm, np.uint8),
                    _copy(lib.dg_aa_nrun_ptr(h, g), n + 1, np.uint64), _copy(lib.dg_aa_nrun_start(h, g), nr, np.uint32),
                    _copy(lib.dg_aa_nrun_end(h, g), nr, np.uint32)))
    return out


def generate(spec: GenSpec, keep_handle: bool = False) -> Dataset:
    lib = _lib()
    h = lib.dg_generate(spec.n_rows, spec.genome_length, spec.seed, int(spec.lineage_model), spec.sub_lambda,
                        spec.n_nodes, spec.zipf_lineage, spec.node_lambda, spec.private_lambda, spec.del_prob,
                        spec.del_max, spec.n_lambda, spec.n_mean, spec.n_countries, spec.zipf_country,
                        spec.n_days, spec.date_base, spec.null_seq_frac, spec.null_meta_frac)
    if not h:
        raise ValueError(f"invalid generator parameters: {spec}")
    n = lib.dg_num_rows(h)
    L = lib.dg_genome_length(h)
    nm = lib.dg_num_mutations(h)
    nr = lib.dg_num_nruns(h)
    nl = lib.dg_num_lineages(h)
    ds = Dataset(
        spec=spec,
        reference=_copy(lib.dg_reference(h), L, np.uint8),
        is_null=_copy(lib.dg_is_null(h), n, np.uint8),
        cov_start=_copy(lib.dg_cov_start(h), n, np.uint32),
        cov_end=_copy(lib.dg_cov_end(h), n, np.uint32),
        mut_ptr=_copy(lib.dg_mut_ptr(h), n + 1, np.uint64),
        mut_pos=_copy(lib.dg_mut_pos(h), nm, np.uint32),
        mut_sym=_copy(lib.dg_mut_sym(h), nm, np.uint8),
        nrun_ptr=_copy(lib.dg_nrun_ptr(h), n + 1, np.uint64),
        nrun_start=_copy(lib.dg_nrun_start(h), nr, np.uint32),
        nrun_end=_copy(lib.dg_nrun_end(h), nr, np.uint32),
        country=_copy(lib.dg_country(h), n, np.uint32),
        country_null=_copy(lib.dg_country_null(h), n, np.uint8),
        date=_copy(lib.dg_date(h), n, np.int32),
        date_null=_copy(lib.dg_date_null(h), n, np.uint8),
        lineage=_copy(lib.dg_lineage(h), n, np.uint32),
        lineage_null=_copy(lib.dg_lineage_null(h), n, np.uint8),
        lineage_parent=_copy(lib.dg_lineage_parent(h), nl, np.int32),
        lineage_weight=_copy(lib.dg_lineage_weight(h), nl, np.float64),
        country_weight=_copy(lib.dg_country_weight(h), spec.n_countries, np.float64),
    )
    cperm = _copy(lib.dg_country_name_perm(h), spec.n_countries, np.uint32)
    lperm = _copy(lib.dg_lineage_name_perm(h), nl, np.uint32)
    ds.country_names = [f"country_{int(cperm[v]):02d}" for v in range(spec.n_countries)]
    ds.lineage_names = [f"L{int(lperm[v]):04d}" for v in range(nl)]
    if keep_handle:
        ds._handle = h
    else:
        lib.dg_free(h)
    return ds


def generate_config(index: int, n_rows: int | None = None, keep_handle: bool = False) -> Dataset:
    spec = CONFIGS[index]
    if n_rows is not None:
        spec = GenSpec(**{**spec.__dict__, "n_rows": n_rows})
    return generate(spec, keep_handle=keep_handle)
