"""Copy the reference's example PLINK trio (data/example/example.{bed,bim,fam}: 150 strains x 1,500 variants)
into a small .npz fixture for the .bed decoder tests.  Data only -- no reference code is imported or run.

Run in the build container only (the reference tree is not present elsewhere):

    python tests/golden/make_bed_fixture.py /root/reference
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def main(ref_root: str) -> None:
    ex = Path(ref_root) / "data" / "example" / "example"
    raw = np.frombuffer(ex.with_suffix(".bed").read_bytes(), dtype=np.uint8)
    fam = [ln.split() for ln in ex.with_suffix(".fam").read_text().splitlines() if ln.strip()]
    bim = [ln.split() for ln in ex.with_suffix(".bim").read_text().splitlines() if ln.strip()]
    np.savez_compressed(
        HERE / "example_bed.npz",
        bed=raw,                                              # the whole file, magic bytes included
        fam=np.array([" ".join(r) for r in fam]),             # .fam lines
        bim=np.array(["\t".join(r) for r in bim]),            # .bim lines
    )
    print(f"wrote example_bed.npz: {raw.size} bed bytes, {len(fam)} strains, {len(bim)} variants")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
