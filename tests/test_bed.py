"""CPU tests of the .bed decoder's oracle and the host-side .bim/.fam handling (no GPU needed)."""
from __future__ import annotations

import numpy as np

from conftest import GOLDEN, golden
from oracle import bed as obed


def test_oracle_decodes_the_format_specification():
    """PLINK 1 SNP-major records, built by hand: codes 00 = hom first allele, 01 = missing, 10 = het,
    11 = hom second allele, individual i at bits 2 (i % 4) of byte i // 4."""
    # 5 individuals, 2 variants.  variant 0 codes [00, 01, 10, 11, 00]; variant 1 codes [11, 11, 10, 00, 01]
    v0 = [0b11100100, 0b00000000]
    v1 = [0b00101111, 0b00000001]
    raw = np.array([0x6C, 0x1B, 0x01] + v0 + v1, dtype=np.uint8)
    Z = obed.decode(raw, 5, 2)
    want = np.array([[2, 0], [np.nan, 0], [1, 1], [0, 2], [2, np.nan]])
    np.testing.assert_array_equal(Z, want)
    Z2 = obed.decode(raw, 5, 2, count_A1=False)
    np.testing.assert_array_equal(Z2, 2.0 - want)


def test_oracle_encode_decode_round_trip():
    rng = np.random.default_rng(7)
    for n in (1, 3, 4, 37, 150):
        Z = rng.integers(0, 3, size=(n, 11)).astype(np.float64)
        Z[rng.random((n, 11)) < 0.1] = np.nan
        for c in (True, False):
            np.testing.assert_array_equal(obed.decode(obed.encode(Z, count_A1=c), n, 11, count_A1=c), Z)


def test_example_fixture_matches_reference_expectations():
    """The reference's example panel (tests/test_io.py:96-122): 150 x 1,500 dosages in {0, 1, 2}, no missing,
    chromosomes 1..16; and the decoded matrix is the panel every other golden vector was computed on."""
    fx = golden("example_bed.npz")
    Z = obed.decode(fx["bed"], len(fx["fam"]), len(fx["bim"]))
    assert Z.shape == (150, 1500)
    assert set(np.unique(Z).tolist()) == {0.0, 1.0, 2.0}
    chrom = np.array([float(ln.split("\t")[0]) for ln in fx["bim"]])
    assert set(chrom.astype(int).tolist()) == set(range(1, 17))
    ex = golden("example_inputs.npz")
    np.testing.assert_array_equal(Z, ex["Z_raw"].astype(np.float64))
    np.testing.assert_array_equal(chrom, ex["chrom"])


def test_bim_fam_parsing(tmp_path):
    from fasterlmm_amd.bed import _read_text
    prefix = tmp_path / "t"
    (tmp_path / "t.fam").write_text("F1 s1 0 0 1 -9\nF2 s2 0 0 2 -9\n")
    (tmp_path / "t.bim").write_text("1\trs1\t0\t100\tA\tG\nX\trs2\t0.5\t200\tC\tT\nchrMT\trs3\t0\t7\tA\tC\n")
    iid, sid, chrom, pos = _read_text(prefix)
    assert iid == ["s1", "s2"] and sid == ["rs1", "rs2", "rs3"]
    np.testing.assert_array_equal(chrom, [1.0, 23.0, 26.0])
    np.testing.assert_array_equal(pos, [100.0, 200.0, 7.0])
