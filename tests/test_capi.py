"""The C-ABI library loads and exports every entry point declared in include/*.h (no GPU work)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> set[str]:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(flmm_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    from fasterlmm_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 9
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding types exactly the declared set
    assert set(_lib.SIGNATURES) == names


def test_abi_version_and_error_channel():
    from fasterlmm_amd import _lib
    lib = _lib.load()
    assert lib.flmm_abi_version() == 3
    # argument validation runs on the host and reports through flmm_last_error without touching a GPU
    rc = lib.flmm_max_reduce(None, 1, 0, 4, None, None)
    assert rc == 1
    assert b"flmm_max_reduce" in lib.flmm_last_error()


def test_bad_arguments_raise_value_error():
    import pytest
    from fasterlmm_amd import _lib
    with pytest.raises(ValueError, match="n_grid"):
        _lib.call("flmm_delta_weights", None, 10, None, 10, 128, 500, -10.0, 10.0, None, None, None)
    with pytest.raises(ValueError, match="precision"):
        _lib.call("flmm_wald_scan", 7, *([None] * 6), 128, *([None] * 6), 128, 128, 1, 1, 0, None, None, 1.0,
                  None, None, None, None, None, None, None, None)
    with pytest.raises(ValueError, match="split-FP16"):
        _lib.call("flmm_pack_planes", 2, 8, 4, 4, 1, 4, None, None, 16, 16, None, 128, 32, None)
