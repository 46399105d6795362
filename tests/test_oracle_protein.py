"""Pins of the oracle's AssortativePair match value (CPU).

* MD5 (orc_md5) against the RFC 1321 test suite and hashlib on random
  messages of every length 0..300 (all padding cases: one block, two blocks,
  exactly 55/56/63/64 bytes).
* The proteome (orc_proteome) against the reference's own ProteinBuilder
  (QHG4/genes/ProteinBuilder.cpp, compiled with the system's OpenSSL MD5 into
  oracle/_ref/libqhg_ref.so by `make -C oracle ref`) on random genomes of
  sizes that cut strands and words at every offset, and on genomes written
  codon by codon for the edge cases: no Met, a Met in the last codon (an
  empty protein closed by the strand's end), Met-stop pairs filling a strand
  (the most proteins a genome can hold), a protein running into the strand's
  end, a partial codon at the end, a Met inside a protein.  The digests are
  compared as multisets (the reference orders them by pointer address).
* The same proteomes from a Python restatement of the translation with
  hashlib's MD5 (no reference needed).
* countProtHashMatches (orc_prot_matches) against brute force, duplicates
  included.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

from oracle.oracle import lib as olib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libqhg_ref.so")

U8P = C.POINTER(C.c_uint8)
U64P = C.POINTER(C.c_uint64)

# the genetic code in the reference's numbering (codon X*16 + Y*4 + Z with
# A=0, C=1, G=2, T=3), written out from the standard code: Met = 10, stop = 20
AMINO = {"Ala": 0, "Cys": 1, "Asp": 2, "Glu": 3, "Phe": 4, "Gly": 5, "His": 6, "Ile": 7, "Lys": 8, "Leu": 9,
         "Met": 10, "Asn": 11, "Pro": 12, "Gln": 13, "Arg": 14, "Ser": 15, "Thr": 16, "Val": 17, "Trp": 18,
         "Tyr": 19, "Stp": 20}
STANDARD_CODE = {
    "TTT": "Phe", "TTC": "Phe", "TTA": "Leu", "TTG": "Leu", "CTT": "Leu", "CTC": "Leu", "CTA": "Leu", "CTG": "Leu",
    "ATT": "Ile", "ATC": "Ile", "ATA": "Ile", "ATG": "Met", "GTT": "Val", "GTC": "Val", "GTA": "Val", "GTG": "Val",
    "TCT": "Ser", "TCC": "Ser", "TCA": "Ser", "TCG": "Ser", "CCT": "Pro", "CCC": "Pro", "CCA": "Pro", "CCG": "Pro",
    "ACT": "Thr", "ACC": "Thr", "ACA": "Thr", "ACG": "Thr", "GCT": "Ala", "GCC": "Ala", "GCA": "Ala", "GCG": "Ala",
    "TAT": "Tyr", "TAC": "Tyr", "TAA": "Stp", "TAG": "Stp", "CAT": "His", "CAC": "His", "CAA": "Gln", "CAG": "Gln",
    "AAT": "Asn", "AAC": "Asn", "AAA": "Lys", "AAG": "Lys", "GAT": "Asp", "GAC": "Asp", "GAA": "Glu", "GAG": "Glu",
    "TGT": "Cys", "TGC": "Cys", "TGA": "Stp", "TGG": "Trp", "CGT": "Arg", "CGC": "Arg", "CGA": "Arg", "CGG": "Arg",
    "AGT": "Ser", "AGC": "Ser", "AGA": "Arg", "AGG": "Arg", "GGT": "Gly", "GGC": "Gly", "GGA": "Gly", "GGG": "Gly",
}
NUC = "ACGT"


def code_of(idx):
    return AMINO[STANDARD_CODE[NUC[idx // 16] + NUC[(idx // 4) % 4] + NUC[idx % 4]]]


def o():
    return olib()


def md5(b):
    arr = (C.c_uint8 * max(1, len(b))).from_buffer_copy(b if b else b"\0")
    out = (C.c_uint8 * 16)()
    o().orc_md5(arr, len(b), out)
    return bytes(out)


def words_of(strands, genome_size):
    """Two strands of nucleotide values (0..3) into genome words (2 bits each,
    from the low end)."""
    nb = (genome_size + 31) // 32
    g = np.zeros(2 * nb, dtype=np.uint64)
    for a, seq in enumerate(strands):
        for i, v in enumerate(seq[:genome_size]):
            g[a * nb + i // 32] |= np.uint64(v) << np.uint64(2 * (i % 32))
    return g


def nucs_of(codons, genome_size, fill=0):
    seq = [NUC.index(ch) for c in codons for ch in c]
    return (seq + [fill] * genome_size)[:genome_size]


def oracle_proteome(g, genome_size):
    mx = 2 * ((genome_size // 3 + 1) // 2 + 1)
    out = (C.c_uint8 * (16 * mx))()
    n = o().orc_proteome(g.ctypes.data_as(U64P), len(g), genome_size, out, mx)
    assert n <= mx
    return sorted(bytes(out[16 * i:16 * i + 16]) for i in range(n))


def python_proteome(g, genome_size):
    """Restatement of translateGenomeHash with hashlib: per strand, codons
    from the first nucleotide; Met opens, stop or the strand's end closes."""
    if genome_size <= 3:
        return []
    nb = len(g) // 2
    out = []
    for a in range(2):
        nuc = [int((int(g[a * nb + i // 32]) >> (2 * (i % 32))) & 3) for i in range(genome_size)]
        cod = [code_of(nuc[3 * k] * 16 + nuc[3 * k + 1] * 4 + nuc[3 * k + 2]) for k in range(genome_size // 3)]
        prot = None
        for aa in cod:
            if prot is None:
                if aa == AMINO["Met"]:
                    prot = []
            elif aa == AMINO["Stp"]:
                out.append(hashlib.md5(bytes(prot)).digest())
                prot = None
            else:
                prot.append(aa)
        if prot is not None:
            out.append(hashlib.md5(bytes(prot)).digest())
    return sorted(out)


def reference_proteome(g, genome_size):
    L = C.CDLL(REF)
    L.ref_protein_hashes.argtypes = [U64P, C.c_int, U8P, C.c_int]
    L.ref_protein_hashes.restype = C.c_int
    mx = 2 * ((genome_size // 3 + 1) // 2 + 1)
    out = (C.c_uint8 * (16 * mx))()
    n = L.ref_protein_hashes(g.ctypes.data_as(U64P), genome_size, out, mx)
    if n < 0:
        return None
    assert n <= mx
    return sorted(bytes(out[16 * i:16 * i + 16]) for i in range(n))


RFC1321 = [
    (b"", "d41d8cd98f00b204e9800998ecf8427e"),
    (b"a", "0cc175b9c0f1b6a831c399e269772661"),
    (b"abc", "900150983cd24fb0d6963f7d28e17f72"),
    (b"message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
    (b"abcdefghijklmnopqrstuvwxyz", "c3fcd3d76192e4007dfb496cca67e13b"),
    (b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789", "d174ab98d277d9f5a5611c2c9f419d9f"),
    (b"1234567890" * 8, "57edf4a22be3c955ac49da2e2107b67a"),
]


@pytest.mark.parametrize("msg,hexd", RFC1321)
def test_md5_rfc1321_suite(msg, hexd):
    assert md5(msg).hex() == hexd


def test_md5_every_length_against_hashlib():
    rng = np.random.default_rng(3)
    for n in range(0, 301):
        b = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        assert md5(b) == hashlib.md5(b).digest(), n


def test_code_table_is_the_standard_code():
    # Met, the three stops and a few anchors of the reference's ordering
    assert code_of(0 * 16 + 3 * 4 + 2) == AMINO["Met"]      # ATG
    assert {code_of(48), code_of(50), code_of(56)} == {AMINO["Stp"]}   # TAA TAG TGA
    assert sum(code_of(i) == AMINO["Stp"] for i in range(64)) == 3


EDGE = {
    "no_met": (["AAA"] * 30, ["CCC"] * 30, 90),
    "met_last_codon": (["AAA"] * 29 + ["ATG"], ["GGG"] * 30, 90),
    "met_stop_pairs": (["ATG", "TAA"] * 15, ["ATG", "TGA"] * 15, 90),
    "into_strand_end": (["AAA", "ATG"] + ["GCT"] * 28, ["ATG", "CAT", "TAG"] + ["AAA"] * 27, 90),
    "partial_codon": (["ATG", "GCT", "GCT"] + ["TTT"] * 29, ["ATG"] * 32, 97),
    "met_inside": (["ATG", "ATG", "ATG", "TAA", "ATG", "TAG"] + ["CCC"] * 25, ["TAA"] * 31, 93),
    "word_straddle": (["GGG"] * 10 + ["ATG"] + ["CAG"] * 9 + ["TGA"] + ["AAA"] * 20, ["ATG"] * 40, 120),
}


@pytest.mark.parametrize("name", list(EDGE))
def test_proteome_edge_cases(name):
    s0, s1, G = EDGE[name]
    g = words_of([nucs_of(s0, G), nucs_of(s1, G)], G)
    want = python_proteome(g, G)
    assert oracle_proteome(g, G) == want
    if name == "met_stop_pairs":
        assert len(want) == 30                      # one per two codons on each strand
    if name == "met_last_codon":
        assert want == [hashlib.md5(b"").digest()]  # closed by the strand's end, empty
    if os.path.exists(REF):
        assert reference_proteome(g, G) == want


@pytest.mark.parametrize("G", [4, 5, 6, 7, 31, 32, 33, 63, 64, 65, 100, 255, 256, 257, 1100])
def test_proteome_random_genomes(G):
    rng = np.random.default_rng(G)
    nb = (G + 31) // 32
    for _ in range(40):
        g = rng.integers(0, 2 ** 63, size=2 * nb, dtype=np.int64).astype(np.uint64)
        want = python_proteome(g, G)
        assert oracle_proteome(g, G) == want
        if os.path.exists(REF):
            assert reference_proteome(g, G) == want


def test_short_genome_has_no_proteome():
    g = words_of([[0, 3, 2], [0, 3, 2]], 3)
    assert oracle_proteome(g, 3) == []
    if os.path.exists(REF):
        assert reference_proteome(g, 3) is None   # ProteinBuilder::init refuses it


def test_prot_matches_counts_every_equal_pair():
    rng = np.random.default_rng(9)
    pool = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(5)]
    for _ in range(200):
        a = [pool[i] for i in rng.integers(0, 5, rng.integers(0, 7))]
        b = [pool[i] for i in rng.integers(0, 5, rng.integers(0, 7))]
        want = sum(x == y for x in a for y in b)
        ab = (C.c_uint8 * max(16, 16 * len(a))).from_buffer_copy(b"".join(a) or bytes(16))
        bb = (C.c_uint8 * max(16, 16 * len(b))).from_buffer_copy(b"".join(b) or bytes(16))
        assert o().orc_prot_matches(ab, len(a), bb, len(b)) == want
