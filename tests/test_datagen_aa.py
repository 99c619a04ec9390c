"""The generator's amino-acid columns: every row's residue diff, re-applied to the translated
reference, equals the direct translation (standard genetic code) of the row's aligned nucleotide
string, with 'X' for codons touching missing data or partly deleted."""
import numpy as np

from silo_amd import datagen
from silo_amd.table import AA_CHARS

CODE = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"
DIGIT = {"T": 0, "C": 1, "A": 2, "G": 3}


def translate_codon(c: str) -> str:
    if c == "---":
        return "-"
    if any(x not in DIGIT for x in c):
        return "X"
    return CODE[16 * DIGIT[c[0]] + 4 * DIGIT[c[1]] + DIGIT[c[2]]]


def test_gene_layout_fits_and_is_ordered():
    lay = datagen.gene_layout(29903)
    assert sum(n for _, _, n in lay) == 9700 and len(lay) == 10
    ends = [s + 3 * n for _, s, n in lay]
    assert all(ends[i] <= lay[i + 1][1] for i in range(len(lay) - 1)) and ends[-1] <= 29903


def test_translation_matches_materialized_rows():
    ds = datagen.generate(datagen.GenSpec(n_rows=400, genome_length=3000, seed=11, lineage_model=True, n_nodes=20,
                                          n_mean=30, del_prob=0.3, null_seq_frac=0.05), keep_handle=True)
    genes = [("g1", 300), ("g2", 450), ("g3", 150)]
    lay = datagen.gene_layout(ds.genome_length, genes, seed=3)
    cols = datagen.translate(ds, lay)
    for (name, start, n), (cname, ref, cs, ce, mptr, mpos, msym, nptr, ns, ne) in zip(lay, cols):
        assert cname == name and len(ref) == n
        refseq = [AA_CHARS[x] for x in ref]
        assert refseq == [translate_codon("".join("-ACGTRYSWKMBDHVN"[b] for b in ds.reference[start + 3 * i:start + 3 * i + 3]))
                          for i in range(n)]
        for r in range(ds.n_rows):
            s = ds.materialize(r)
            want = [translate_codon(s[start + 3 * i:start + 3 * i + 3]) for i in range(n)]
            if ds.is_null[r]:
                continue
            got = ["X"] * n
            got[cs[r]:ce[r]] = refseq[cs[r]:ce[r]]
            for k in range(nptr[r], nptr[r + 1]):
                got[ns[k]:ne[k]] = ["X"] * (ne[k] - ns[k])
            for k in range(mptr[r], mptr[r + 1]):
                assert mpos[k] >= cs[r] and mpos[k] < ce[r] and AA_CHARS[msym[k]] != refseq[mpos[k]]
                got[mpos[k]] = AA_CHARS[msym[k]]
            assert got == want, (name, r)
