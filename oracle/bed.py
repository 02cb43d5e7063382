"""CPU oracle for the PLINK 1 .bed decoder -- TEST INFRASTRUCTURE ONLY.

Restates what the reference reads through pysnptools (fasterlmm/io.py:54-67 ``read_plink`` with
``Bed(count_A1=True)``, fasterlmm/io_stream.py:60-71 ``read_block``): a SNP-major .bed is 3 magic bytes
(0x6C 0x1B 0x01) and one record of ceil(N / 4) bytes per variant, individual i in byte i // 4 at bits
2 (i % 4) .. 2 (i % 4) + 1.  PLINK's codes are 00 = homozygous first (.bim A1) allele, 01 = missing,
10 = heterozygous, 11 = homozygous second allele; count_A1 counts the first allele (00 -> 2, 10 -> 1,
11 -> 0), otherwise the second.  Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.

Pinning: hand-built records from the format specification (tests/test_bed.py), an independent encoder
round trip, and the reference's own expectations for its example panel (tests/test_io.py:96-122: 150 x 1,500
dosages in {0, 1, 2}, chromosomes 1..16).
"""
from __future__ import annotations

import numpy as np

MAGIC = bytes([0x6C, 0x1B, 0x01])


def decode(bed: np.ndarray, n: int, m: int, *, count_A1: bool = True) -> np.ndarray:
    """(n, m) float64 dosages with NaN for missing from the raw file bytes (magic included)."""
    raw = np.asarray(bed, dtype=np.uint8)
    if raw[:3].tobytes() != MAGIC:
        raise ValueError("not a SNP-major PLINK 1 .bed")
    rec = (n + 3) // 4
    body = raw[3:3 + rec * m].reshape(m, rec)
    codes = np.stack([(body >> (2 * k)) & 3 for k in range(4)], axis=2).reshape(m, rec * 4)[:, :n]
    lut = np.array([2.0, np.nan, 1.0, 0.0]) if count_A1 else np.array([0.0, np.nan, 1.0, 2.0])
    return lut[codes].T.copy()


def encode(Z: np.ndarray, *, count_A1: bool = True) -> np.ndarray:
    """Inverse of decode: (n, m) dosages in {0, 1, 2, NaN} -> file bytes (magic included)."""
    Z = np.asarray(Z, dtype=np.float64)
    n, m = Z.shape
    a1 = Z if count_A1 else 2.0 - Z
    code = np.full((m, n), 1, dtype=np.uint8)  # missing
    zt = a1.T
    code[zt == 2.0] = 0
    code[zt == 1.0] = 2
    code[zt == 0.0] = 3
    rec = (n + 3) // 4
    padded = np.zeros((m, rec * 4), dtype=np.uint8)
    padded[:, :n] = code
    q = padded.reshape(m, rec, 4)
    body = (q[..., 0] | (q[..., 1] << 2) | (q[..., 2] << 4) | (q[..., 3] << 6)).astype(np.uint8)
    return np.concatenate([np.frombuffer(MAGIC, dtype=np.uint8), body.reshape(-1)])
