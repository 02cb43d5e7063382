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
    assert lib.flmm_abi_version() == 18
    # argument validation runs on the host and reports through flmm_last_error without touching a GPU
    rc = lib.flmm_max_reduce(None, 1, 0, 4, None, None)
    assert rc == 1
    assert b"flmm_max_reduce" in lib.flmm_last_error()


def test_bad_arguments_raise_value_error():
    import pytest
    from fasterlmm_amd import _lib
    with pytest.raises(ValueError, match="n_grid"):
        _lib.call("flmm_delta_weights", None, 10, None, 10, 128, 500, -10.0, 10.0, None, None, None, None)
    with pytest.raises(ValueError, match="precision"):
        _lib.call("flmm_wald_scan", 7, *([None] * 6), 128, *([None] * 6), 128, 128, 1, 1, 0, None, None, 1.0,
                  None, None, None, None, None, None, None, None)
    with pytest.raises(ValueError, match="split-FP16"):
        _lib.call("flmm_pack_planes", 2, 8, 4, 4, 1, 4, None, None, 16, 16, None, 128, 32, None)
    with pytest.raises(ValueError, match="rows % 64"):
        _lib.call("flmm_gemm_f64", 16, 100, 32, 16, 128, 16, None)
    with pytest.raises(ValueError, match="null pointer"):
        _lib.call("flmm_gemm_f64", None, 64, 32, 16, 128, 16, None)
    with pytest.raises(ValueError, match="neff <= kpad"):
        _lib.call("flmm_rotate_f64", 16, 10, None, 5, 10, 16, 200, 200, 16, 128, 128, None)
    with pytest.raises(ValueError, match="null pointer"):
        _lib.call("flmm_rotate_f64", 16, 10, None, 5, 10, None, 200, 100, 16, 128, 128, None)
    with pytest.raises(ValueError, match="ldz >= m"):
        _lib.call("flmm_gram_f64", 16, 99, 10, 100, 1, 16, 16, 10, None)
    with pytest.raises(ValueError, match="null pointer"):
        _lib.call("flmm_gram_f64", 16, 100, 10, 100, 1, None, 16, 10, None)
    with pytest.raises(ValueError, match="bad sizes"):
        _lib.call("flmm_rescore_f64", 16, 16, 128, 16, 200, 16, 16, 16, 16, 0, 5, None, 10.0, 16, 16, 16, 16, 4, None)
    with pytest.raises(ValueError, match="null pointer"):
        _lib.call("flmm_rescore_f64", 16, None, 128, 16, 100, 16, 16, 16, 16, 0, 5, None, 10.0, 16, 16, 16, 16, 4, None)
    with pytest.raises(ValueError, match="R <= rpad"):
        _lib.call("flmm_factored_weight_planes", 2, 10, 256, 128, 16, 16, 129, 16, 16, 16, 16, 256, None, None)
    with pytest.raises(ValueError, match="q2_scale"):
        _lib.call("flmm_factored_weight_planes", 2, 10, 256, 128, 16, 16, 100, 16, 16, 16, None, 256, None, None)


def test_screen_layout_buckets():
    """The bucketed screen's layout on the host: buckets of width <= 1/5 over the log-delta range (at most 120, else
    no layout), Q-plane rows covering every padded slot, and the checks flmm_delta_refine makes on them."""
    import pytest
    from fasterlmm_amd import _lib, engine
    lib = _lib.load()
    P, p_pad, hf = 1000, 1024, 256
    lay = engine.screen_layout(hf, P, p_pad, (-10.0, 10.0), "cpu")
    assert (lay.nb, lay.t_lo, lay.t_h) == (100, -10.0, 0.2)
    assert lay.q_rows == lib.flmm_screen_q_rows(p_pad, hf, P) == (hf + (P - hf) + 32 * 120 + 255) // 256 * 256
    assert lay.ws.numel() == lib.flmm_screen_order_bytes(P - hf)
    lay = engine.screen_layout(hf, P, p_pad, (0.1, 0.7), "cpu")
    assert lay.nb == 3 and abs(lay.t_h - 0.2) < 1e-15 and lay.t_h <= 0.2
    assert engine.screen_layout(hf, P, p_pad, (3.0, 3.0), "cpu").nb == 1
    assert engine.screen_layout(hf, P, p_pad, (-12.0, 12.0), "cpu").nb == 120
    assert engine.screen_layout(hf, P, p_pad, (-12.0, 12.01), "cpu") is None  # 121 buckets
    assert engine.screen_layout(None, P, p_pad, (-10.0, 10.0), "cpu") is None
    assert engine.screen_layout(P, P, p_pad, (-10.0, 10.0), "cpu") is None
    with engine.configure(screen="general"):
        assert engine.screen_layout(hf, P, p_pad, (-10.0, 10.0), "cpu") is None
    assert lib.flmm_screen_q_rows(p_pad, P, P) == p_pad and lib.flmm_screen_q_rows(p_pad, 0, p_pad + 1) == -1
    assert lib.flmm_screen_work_bytes(256, 1024) == 8 * 256 * (1024 + 256)
    # flmm_delta_refine's layout checks (host only: they fail before any launch)
    args = [2, 0, 16, P, p_pad, 128, 16, 100, 16, 100, 900.0, 64, -10.0, 10.0, 16, 16, 16, 16, 16, 16, 16, 16, 16,
            16, 16, hf]
    with pytest.raises(ValueError, match="q_rows below"):
        _lib.call("flmm_delta_refine", *args, p_pad, -10.0, 0.2, 100, 16, None)
    q = lib.flmm_screen_q_rows(p_pad, hf, P)
    with pytest.raises(ValueError, match="t_h <= 1/5"):
        _lib.call("flmm_delta_refine", *args, q, -10.0, 0.25, 80, 16, None)
    with pytest.raises(ValueError, match="t_h <= 1/5"):
        _lib.call("flmm_delta_refine", *args, q, -10.0, 0.2, 121, 16, None)
    with pytest.raises(ValueError, match="multiple of 256"):
        _lib.call("flmm_delta_refine", *(args[:-1] + [128]), q, -10.0, 0.2, 100, 16, None)


def _fake_group(_lib, mc=300, P=200, p_pad=256, n_real=3, base=0x10000):
    """A scan group whose pointers are distinct, 16-byte aligned, never dereferenced addresses."""
    g = _lib.ScanGroup()
    for i, (name, typ) in enumerate(g._fields_):
        if typ is ctypes.c_void_p:
            setattr(g, name, base + 256 * i)
    g.mc_pad = (mc + 127) // 128 * 128
    g.p_pad, g.kpad, g.mc, g.P, g.n_real, g.nu = p_pad, 1024, mc, P, n_real, 990.0
    return g


def test_scan_table_layout_and_validation():
    """flmm_scan_table_fill is host-only: header, tile prefix sums and the packed group records."""
    import numpy as np
    import pytest
    from fasterlmm_amd import _lib
    lib = _lib.load()
    assert ctypes.sizeof(_lib.ScanGroup) == 30 * 8
    groups = (_lib.ScanGroup * 3)(_fake_group(_lib, 300, 200), _fake_group(_lib, 128, 256, base=0x20000),
                                  _fake_group(_lib, 0, 256, base=0x30000))
    nbytes = lib.flmm_scan_table_bytes(3)
    assert nbytes == (4 + 4) * 8 + 3 * 30 * 8
    for variant, tp in ((_lib.SCAN_SQUARE, 128), (_lib.SCAN_WIDE, 256)):
        buf = np.zeros(nbytes // 8, np.int64)
        _lib.call("flmm_scan_table_fill", _lib.PREC_SPLIT_FP16, variant, ctypes.addressof(groups), 3,
                  buf.ctypes.data)
        tiles = [3 * -(-200 // tp), 1 * -(-256 // tp), 0]
        assert list(buf[:4]) == [3, sum(tiles), variant, 0]
        assert list(buf[4:8]) == [0, tiles[0], tiles[0] + tiles[1], sum(tiles)]
        assert buf[8:].tobytes() == bytes(groups)
    bad = (_lib.ScanGroup * 1)(_fake_group(_lib, p_pad=384))
    with pytest.raises(ValueError, match="p_pad % 256"):
        _lib.call("flmm_scan_table_fill", 0, _lib.SCAN_WIDE, ctypes.addressof(bad), 1, buf.ctypes.data)
    _lib.call("flmm_scan_table_fill", 0, _lib.SCAN_SQUARE, ctypes.addressof(bad), 1, buf.ctypes.data)
    bad[0].q2_lo += 8
    with pytest.raises(ValueError, match="group 0: operand planes must be 16-byte aligned"):
        _lib.call("flmm_scan_table_fill", 0, _lib.SCAN_SQUARE, ctypes.addressof(bad), 1, buf.ctypes.data)
    # factored weights (kpad2 > 0) run on the wide variant only, with kpad2 a multiple of 128 below kpad
    fac = (_lib.ScanGroup * 1)(_fake_group(_lib))
    with pytest.raises(ValueError, match="factored variant needs kpad2"):
        _lib.call("flmm_scan_table_fill", 0, _lib.SCAN_WIDE_FACTORED, ctypes.addressof(fac), 1, buf.ctypes.data)
    fac[0].kpad2 = 128
    _lib.call("flmm_scan_table_fill", 0, _lib.SCAN_WIDE_FACTORED, ctypes.addressof(fac), 1, buf.ctypes.data)
    assert list(buf[:3]) == [1, 3 * 1, _lib.SCAN_WIDE_FACTORED]
    for v in (_lib.SCAN_WIDE, _lib.SCAN_SQUARE):
        with pytest.raises(ValueError, match="kpad2 > 0 needs the factored variant"):
            _lib.call("flmm_scan_table_fill", 0, v, ctypes.addressof(fac), 1, buf.ctypes.data)
    fac[0].kpad2 = 1024
    with pytest.raises(ValueError, match="factored variant needs kpad2"):
        _lib.call("flmm_scan_table_fill", 0, _lib.SCAN_WIDE_FACTORED, ctypes.addressof(fac), 1, buf.ctypes.data)
    with pytest.raises(ValueError, match="blocks must be a positive multiple of 8"):
        _lib.call("flmm_wald_scan_grouped", 0, 1, buf.ctypes.data, 100, None)


def test_kinship_setting_is_validated_before_any_launch():
    """engine.gram rejects an unknown kinship setting, and the split Gram outside the split precisions, on the host
    (no device memory is touched)."""
    import pytest
    import torch
    from fasterlmm_amd import engine
    Z = torch.zeros(3, 5, dtype=torch.float64)
    with engine.configure(kinship="bogus"):
        with pytest.raises(ValueError, match="kinship"):
            engine.gram(Z)
    with engine.configure(kinship="split", precision="fp32"):
        with pytest.raises(ValueError, match="split precision"):
            engine.gram(Z)
